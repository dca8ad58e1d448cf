"""GPU parity: the HIP path, called through the C-ABI, against the oracle and
the golden fixtures.  U1 and the tie term must be bit-identical; asymptotic
and exact p within 1e-12 relative of the FP64 oracle; group means within
1e-6 relative."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from fixtures import group_mean_cases, mwu_cases, rel_close, small_config_expected
from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu

ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}
ALT = {"less": 0, "greater": 1, "two-sided": 2}
METH = {"exact": 1, "asymptotic": 2}
P_RTOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_stats(w, mode, alt="two-sided", method="asymptotic", tie=True, cc=True):
    opts = _capi.make_opts(tie_correction=tie, use_continuity=cc, zero_handling=ZH[mode],
                           alternative=ALT[alt], method=METH[method])
    return _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                         opts, n_rows=w["n_rows"], with_stats=True)


ORACLE_THREADS = min(16, os.cpu_count() or 1)


def oracle_stats(w, mode, alt="two-sided", method="asymptotic", tie=True, cc=True):
    return oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                      n_rows=w["n_rows"], tie_correction=tie, use_continuity=cc,
                      zero_handling=mode, alternative=alt, method=method, with_stats=True,
                      threads=ORACLE_THREADS)


def assert_parity(g, o, p_rtol=P_RTOL):
    U1g, Pg, Tg, n1g, n2g = g
    U1o, Po, To, n1o, n2o = o
    assert np.array_equal(n1g, n1o) and np.array_equal(n2g, n2o), "sample sizes differ"
    bad = np.nonzero(U1g != U1o)[0]
    assert len(bad) == 0, f"U1 differs at {bad[:5]}: {U1g[bad[:5]]} vs {U1o[bad[:5]]}"
    bad = np.nonzero(Tg != To)[0]
    assert len(bad) == 0, f"tie differs at {bad[:5]}: {Tg[bad[:5]]} vs {To[bad[:5]]}"
    assert rel_close(Pg, Po, rtol=p_rtol), float(np.max(np.abs(Pg - Po) / np.maximum(Po, 1e-300)))


CASES = mwu_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_cases(case):
    say(f"golden {case['name']}")
    w = dict(data=case["data"], indices=case["indices"], indptr=case["indptr"],
             group_id=case["group_id"], n_targets=case["n_targets"], n_rows=case["n_rows"])
    g = gpu_stats(w, case["zero_handling"], case["alternative"], case["method"],
                  case["tie_correction"], case["use_continuity"])
    assert np.array_equal(g[0], case["U1"])
    assert np.array_equal(g[2], case["tie"])
    assert rel_close(g[1], case["P"], rtol=1e-12)
    o = oracle_stats(w, case["zero_handling"], case["alternative"], case["method"],
                     case["tie_correction"], case["use_continuity"])
    assert_parity(g, o)


def test_small_config_and_wrapper():
    say("small config")
    w = synthetic.small_case()
    exp = small_config_expected()
    g = gpu_stats(w, "min")
    assert np.array_equal(g[0], exp["U1"]) and np.array_equal(g[2], exp["tie"])
    assert_parity(g, oracle_stats(w, "min"))
    from hpdex_amd.backend import group_mean, mannwhitneyu
    M = sp.csc_matrix((w["data"], w["indices"], w["indptr"]), shape=(w["n_rows"], w["n_cols"]))
    U1, P = mannwhitneyu(M, w["group_id"], 4, use_continuity=True, tie_correction=True,
                         zero_handling="min")
    assert not U1.flags.writeable and not P.flags.writeable and U1.dtype == np.float64
    assert np.array_equal(U1, exp["U1"]) and rel_close(P, exp["P"])
    gm = group_mean(M, w["group_id"], 5, include_zeros=True)
    assert rel_close(gm, exp["group_mean"], rtol=1e-6, atol=0)


def test_kernel_module_default_arguments():
    """kernel.mannwhitneyu with every default of the extension module
    (module.cpp:215-228: tie correction on, continuity off, fast_norm=True,
    zero_handling none, two-sided, asymptotic): fast_norm runs the precise
    erfc path, so p matches the oracle's precise p to 1e-12."""
    say("kernel module defaults")
    from hpdex_amd.backend.bin import kernel
    w = synthetic.small_case()
    U1, P = kernel.mannwhitneyu(w["data"], w["indices"], w["indptr"], w["group_id"], 4)
    oU1, oP = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], 4,
                         tie_correction=True, use_continuity=False, zero_handling="none")
    assert np.array_equal(U1, oU1)
    assert rel_close(P, oP, rtol=P_RTOL)
    U1f, Pf = kernel.mannwhitneyu(w["data"], w["indices"], w["indptr"], w["group_id"], 4,
                                  fast_norm=False)
    assert np.array_equal(U1f, U1) and np.array_equal(Pf, P)


NEGZERO = [c for c in CASES if c["name"].startswith("negzero")]


@pytest.mark.parametrize("case", NEGZERO, ids=[c["name"] for c in NEGZERO])
def test_negative_zero_float32_paths(case, restore_tuning):
    """Stored -0.0 beside +0.0 and negative values, as float32: integer
    columns on the count path (packed entries and rows + values) and on the
    sort path's histogram branch, continuous columns on the sort path; every
    route equals the oracle and the scipy-derived golden expectation."""
    say(f"negative zero {case['name']}")
    d32 = case["data"].astype(np.float32)
    assert np.any((d32 == 0) & np.signbit(d32))
    w = dict(data=d32, indices=case["indices"], indptr=case["indptr"],
             group_id=case["group_id"], n_targets=case["n_targets"], n_rows=case["n_rows"])
    mode = case["zero_handling"]
    o = oracle_stats(w, mode)
    assert np.array_equal(o[0], case["U1"]) and np.array_equal(o[2], case["tie"])
    routes = [(1, 40448), (0, 40448), (1, 0)] if case["name"].endswith("int") else [(1, 40448)]
    for packed, lds in routes:
        _capi.set_tuning(count_packed=packed, count_lds_bytes=lds)
        m, p = _plan_of(w, mode)
        p.run(1)
        info = p.info()
        if case["name"].endswith("int"):
            assert info["count_columns"] == (len(case["indptr"]) - 1 if lds else 0)
        else:
            assert info["split_columns"] == len(case["indptr"]) - 1
        assert_parity(p.results(with_stats=True), o)
        p.close(); m.close()
    assert_parity(gpu_stats(w, mode), o)


@pytest.mark.parametrize("k", range(4))
def test_group_mean_golden(k):
    say(f"group_mean {k}")
    c = group_mean_cases()[k]
    out = _capi.group_mean_csc(c["data"], c["indices"], c["indptr"], c["group_id"],
                               c["n_groups"], include_zeros=c["include_zeros"])
    assert rel_close(out, c["expected"], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32,
                                   np.int64, np.uint64, np.float32, np.float64])
def test_all_dtypes(dtype):
    say(f"dtype {np.dtype(dtype).name}")
    rng = np.random.default_rng(11)
    R, C, T = 300, 40, 3
    dense = rng.integers(0, 9, (R, C)) * (rng.random((R, C)) < 0.4)
    if np.dtype(dtype).kind == "i":
        dense = dense - 3 * (rng.random((R, C)) < 0.2)
    M = sp.csc_matrix(dense.astype(dtype))
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    w = dict(data=M.data, indices=M.indices, indptr=M.indptr, group_id=gid, n_targets=T,
             n_rows=R)
    for mode in ("min", "mix", "none"):
        assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))
    gm = _capi.group_mean_csc(M.data, M.indices, M.indptr, gid, T + 1)
    ref = oracle.group_mean(M.data, M.indices, M.indptr, gid, T + 1)
    assert rel_close(gm, ref, rtol=1e-6, atol=1e-12)


def test_error_messages_match_oracle():
    say("errors")
    cases = []
    gid = np.array([0, 1, 1, 1], dtype=np.int32)
    M = sp.csc_matrix(np.arange(1, 9, dtype=np.float32).reshape(4, 2))
    cases.append((M, gid, 1, "min"))
    gid = np.array([0, 0, 0, 1, 1, 1, 2], dtype=np.int32)
    M = sp.csc_matrix(np.arange(1, 15, dtype=np.float32).reshape(7, 2))
    cases.append((M, gid, 2, "mix"))
    dense = np.array([[1, 1], [2, 2], [3, 0], [4, 5], [5, 0]], dtype=np.float32)
    cases.append((sp.csc_matrix(dense), np.array([0, 0, 1, 1, 1], dtype=np.int32), 1, "none"))
    dense = np.array([[1, 0], [0, 2], [3, 1], [4, 5], [5, 6]], dtype=np.float32)
    cases.append((sp.csc_matrix(dense), np.array([0, 0, 1, 1, 1], dtype=np.int32), 1, "none"))
    for M, gid, T, mode in cases:
        with pytest.raises(oracle.OracleError) as eo:
            oracle.mwu(M.data, M.indices, M.indptr, gid, T, zero_handling=mode)
        with pytest.raises(_capi.HpdexError) as eg:
            _capi.mwu_csc(M.data, M.indices, M.indptr, gid, T,
                          _capi.make_opts(zero_handling=ZH[mode]))
        assert str(eg.value) == str(eo.value)
    with pytest.raises(_capi.HpdexError, match="n_targets must be >= 1"):
        _capi.mwu_csc(M.data, M.indices, M.indptr, gid, 0, _capi.make_opts())


def _random_workload(seed, R, C, T, dens, values, ref_frac=0.3):
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, dens))
    data = values(rng, len(rows)).astype(np.float32)
    gid = np.full(R, -1, dtype=np.int32)
    perm = rng.permutation(R)
    nref = int(R * ref_frac)
    gid[perm[:nref]] = 0
    gid[perm[nref:]] = rng.integers(1, T + 1, R - nref)
    return dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R,
                n_cols=C)


@pytest.mark.parametrize("mode", ["min", "max", "mix", "none"])
def test_float_sort_path(mode):
    """Continuous values with negatives: the sort path, large reference slices
    (multi-chunk LDS sort) and target slices above 1024 (chunked ties)."""
    say(f"float sort path {mode}")
    w = _random_workload(5, 12000, 24, 3, 0.5,
                         lambda r, n: np.round(r.normal(0, 2, n), 2))
    assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))


def test_split_scatter_beyond_lds():
    """Columns of ~50,000 entries scattered by group offsets (their groups hold
    more rows than they store): 200 KB of keys do not fit a workgroup's LDS,
    so the scatter writes to HBM; a 5,000-entry column stages in LDS."""
    say("split scatter beyond lds")
    rng = np.random.default_rng(21)
    R, T = 200_000, 30
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.25, 0.25, 0.025]))
    data = np.round(rng.normal(0, 1.5, len(rows)), 2).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:25000]] = 0       # ~6,250 stored reference values per wide column
    gid[perm[25000:150000]] = rng.integers(1, T + 1, 125000)
    w = dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)
    for mode in ("min", "none"):
        assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))


@pytest.mark.parametrize("grouped", [20000, 45000])
def test_dense_split_slots(grouped):
    """Dense columns split in one pass into slots: 20,000 grouped rows keep
    their keys in LDS and write them out coalesced, 45,000 (180 KB of keys)
    write each key to its slot in HBM; explicit zeros leave slot gaps."""
    say(f"dense split slots {grouped}")
    rng = np.random.default_rng(20)
    R, C, T = 46000, 6, 40
    X = np.round(rng.normal(0, 1.5, (C, R)), 1).astype(np.float32)
    X[rng.random((C, R)) < 0.05] = 0.0
    indices = np.tile(np.arange(R, dtype=np.int64), C)
    indptr = np.arange(C + 1, dtype=np.int64) * R
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:6000]] = 0        # within the 8,192 reference keys kept on chip
    gid[perm[6000:grouped]] = rng.integers(1, T + 1, grouped - 6000)
    w = dict(data=X.reshape(-1), indices=indices, indptr=indptr, group_id=gid, n_targets=T,
             n_rows=R)
    for mode in ("min", "none"):
        assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))


@pytest.mark.parametrize("f64,n_ref", [(False, 60000), (True, 20000)], ids=["f32-60k", "f64-20k"])
def test_reference_beyond_chip(f64, n_ref):
    """Reference slices of up to 60,000 four-byte or 20,000 eight-byte keys
    per column, beyond the on-chip sort (8,192 / 4,096 keys): sorted in HBM
    (LDS-sorted runs and merge passes through the job's spill space) and
    ranked from HBM, beside columns whose reference fits the chip; ties
    inside the reference and against targets, negatives; every zero mode;
    plain and batched plans bit-identical to the one-shot call."""
    say(f"reference beyond chip f64={f64} n_ref={n_ref}")
    rng = np.random.default_rng(23)
    R, T = n_ref + 24000, 40
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([1.0, 0.97, 0.6, 0.25, 0.1]))
    data = np.round(rng.normal(0, 3, len(rows)), 2).astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:n_ref]] = 0
    gid[perm[n_ref:]] = rng.integers(1, T + 1, R - n_ref)
    w = dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)
    for mode in ("min", "max", "mix", "none"):
        o = oracle_stats(w, mode)
        g = gpu_stats(w, mode)
        assert_parity(g, o)
        if mode == "min":
            for form in (0, 1):
                m, p = _plan_of(w, mode, form=form)
                p.run(2)
                assert p.info()["split_columns"] == 5
                for a, b in zip(p.results(with_stats=True), g):
                    assert np.array_equal(a, b)
                p.close(); m.close()


@pytest.mark.parametrize("f64", [False, True], ids=["f32", "f64"])
def test_reference_cutover_tuning(f64, restore_tuning):
    """The LDS/global sort cutover as a tuning (ref_lds_keys 8192, 1024, 64):
    reference slices above it are sorted in HBM and ranked through L2, the
    rest on chip; results bit-identical across cutovers and equal to the
    oracle, the tuning refused outside powers of two in [64, 8192]."""
    say(f"reference cutover tuning f64={f64}")
    rng = np.random.default_rng(43)
    R, T = 12000, 30
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([1.0, 0.7, 0.3, 0.1, 0.02]))
    data = np.round(rng.normal(0, 2, len(rows)), 2).astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:2500]] = 0
    gid[perm[2500:]] = rng.integers(1, T + 1, R - 2500)
    w = dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)
    o = oracle_stats(w, "min")
    first = None
    for cap in (8192, 1024, 64):
        _capi.set_tuning(ref_lds_keys=cap)
        g = gpu_stats(w, "min")
        assert_parity(g, o)
        if first is None:
            first = g
        for a, b in zip(g, first):
            assert np.array_equal(a, b)
    for bad in (0, 32, 100, 16384):
        with pytest.raises(_capi.HpdexError, match="ref_lds_keys"):
            _capi.set_tuning(ref_lds_keys=bad)


@pytest.mark.parametrize("mode", ["min", "none"])
def test_float_windows_large_reference(mode):
    """Small targets against a reference of ~5,800 stored values per column:
    a window's sums may pass 2^32, so the target windows scan in 64 bits
    (smaller references take 32-bit scans)."""
    say(f"float windows large reference {mode}")
    rng = np.random.default_rng(19)
    R, C, T = 9000, 8, 120
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, 0.97))
    data = np.round(rng.normal(0, 1.0, len(rows)), 1).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:6000]] = 0
    at = 6000
    for g, n in enumerate(rng.integers(3, 31, T), start=1):
        gid[perm[at:at + n]] = g
        at += n
    w = dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)
    assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))


@pytest.mark.parametrize("mode", ["min", "mix", "none"])
def test_integer_hist_path(mode):
    say(f"integer hist path {mode}")
    w = _random_workload(6, 20000, 30, 25, 0.3,
                         lambda r, n: r.negative_binomial(0.8, 0.2, n) - (mode != "min") * 3)
    assert_parity(gpu_stats(w, mode), oracle_stats(w, mode))


def test_wide_integer_range_uses_sort_path():
    say("wide integer range")
    w = _random_workload(7, 5000, 20, 6, 0.4, lambda r, n: r.integers(-100000, 100000, n))
    assert_parity(gpu_stats(w, "mix"), oracle_stats(w, "mix"))


@pytest.mark.parametrize("mode", ["min", "none"])
@pytest.mark.parametrize("alt", ["two-sided", "less", "greater"])
def test_exact_path(mode, alt):
    say(f"exact {mode} {alt}")
    w = synthetic.exact_case(n_cells=3000, n_genes=40, n_ref=60, n_targets=60)
    g = gpu_stats(w, mode, alt, "exact", True, False)
    o = oracle_stats(w, mode, alt, "exact", True, False)
    assert_parity(g, o)


@pytest.fixture
def restore_tuning():
    t = _capi.get_tuning()
    yield
    _capi.set_tuning(t.hist_max_bins, t.size_ordered, t.count_lds_bytes, t.count_block,
                     t.count_packed, t.split_arrival, t.ref_lds_keys)


def _plan_of(w, mode, form=1, stats=True):
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=ZH[mode])
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    return m, _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=form, stats=stats)


@pytest.mark.parametrize("lds", [0, 16 * 1024, 24 * 1024, 152 * 1024])
def test_count_kernel_lds_sizes(lds, restore_tuning):
    """The one-pass count kernel at several LDS sizes: 0 sends every column to
    the split path; 16 and 24 KiB split the 40 targets of the wider columns
    into slices; 152 KiB takes each column in one workgroup."""
    w = _random_workload(11, 9000, 16, 40, 0.6,
                         lambda r, n: r.negative_binomial(0.8, 0.02, n) - 2)
    for packed in (0, 1):
        # read as rows and values the bins are the value range; packed, the
        # distinct values (fewer, so at most as many slices)
        _capi.set_tuning(count_lds_bytes=lds, count_packed=packed)
        for mode in ("min", "max", "mix", "none"):
            say(f"count kernel lds={lds} packed={packed} mode={mode}")
            m, p = _plan_of(w, mode)
            p.run(1)
            info = p.info()
            assert info["count_columns"] == (16 if lds else 0)
            assert info["count_packed_columns"] == (16 if lds and packed else 0)
            assert info["handed_over"] == 0
            if lds == 152 * 1024:
                assert info["count_blocks"] == 16
            elif lds and not packed:
                assert info["count_blocks"] > 16
            assert_parity(p.results(with_stats=True), oracle_stats(w, mode))
            p.close(); m.close()


def test_count_kernel_packed_entries(restore_tuning):
    """The count kernel reading a column's packed entries (one word per entry:
    row and value code) and reading its rows and values give the same results,
    equal to the oracle's: explicit zeros, NaN and +-Inf (the skip code),
    rows outside [0, R), at R = 300,000 (13 code bits), where a column of
    8,190 bins is packed and one of 8,191 is not."""
    rng = np.random.default_rng(17)
    R, T = 300_000, 5
    rows, ip = synthetic._bernoulli_rows(rng, R, np.array([0.03, 0.03, 0.05, 0.05, 0.2]))
    rows = rows.copy()
    d = rng.integers(1, 8190, len(rows)).astype(np.float32)
    d[ip[0]], d[ip[0] + 1] = 1.0, 8190.0
    d[ip[1]], d[ip[1] + 1] = 1.0, 8191.0
    d[ip[2]:] = rng.negative_binomial(0.8, 0.3, len(d) - ip[2]) - 1
    d[ip[2]:ip[3]:13] = np.nan
    d[ip[3]:ip[4]:17] = np.inf
    d[ip[3] + 5] = -np.inf
    rows[ip[4] + 3] = -4
    rows[ip[4] + 11] = R + 9
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    w = dict(data=d, indices=rows, indptr=ip, group_id=gid, n_targets=T, n_rows=R, n_cols=5)
    for mode in ("min", "none", "mix"):
        say(f"count kernel packed entries mode={mode}")
        o = oracle_stats(w, mode)
        res = {}
        for packed in (1, 0):
            # 152 KiB: the 8,190- and 8,191-bin columns fit (two target slices)
            _capi.set_tuning(count_packed=packed, count_lds_bytes=152 * 1024)
            m, p = _plan_of(w, mode)
            p.run(2)
            info = p.info()
            assert info["count_columns"] == 5 and info["handed_over"] == 0
            assert info["count_packed_columns"] == (4 if packed else 0)
            res[packed] = p.results(with_stats=True)
            p.close(); m.close()
            assert_parity(res[packed], o)
        assert all(np.array_equal(a, b) for a, b in zip(res[0], res[1]))


def test_count_kernel_rows_beyond_24_bits():
    """R = 2^24 + 100 rows: the label positions take the 32-bit multiply; 25
    row bits leave 7 code bits, so a 100-bin column is packed and a 300-bin
    column is read as rows and values; rows up to R - 1."""
    say("count kernel rows beyond 2^24")
    rng = np.random.default_rng(18)
    R, T, n = (1 << 24) + 100, 4, 6000
    rows0 = np.sort(rng.choice(R, n, replace=False))
    rows0[-3:] = [R - 3, R - 2, R - 1]
    rows1 = np.sort(rng.choice(R, n, replace=False))
    rows = np.concatenate([rows0, rows1]).astype(np.int64)
    ip = np.array([0, n, 2 * n], dtype=np.int64)
    d = np.concatenate([rng.integers(1, 101, n), rng.integers(1, 301, n)]).astype(np.float32)
    d[0], d[1] = 1.0, 100.0
    d[n], d[n + 1] = 1.0, 300.0
    gid = np.full(R, -1, np.int32)
    pick = rng.choice(R, 400_000, replace=False)
    gid[pick] = rng.integers(0, T + 1, len(pick))
    gid[rows] = rng.integers(0, T + 1, len(rows))
    w = dict(data=d, indices=rows, indptr=ip, group_id=gid, n_targets=T, n_rows=R, n_cols=2)
    for mode in ("min", "none"):
        m, p = _plan_of(w, mode)
        p.run(1)
        info = p.info()
        assert info["count_columns"] == 2 and info["count_packed_columns"] == 1
        assert info["handed_over"] == 0
        assert_parity(p.results(with_stats=True), oracle_stats(w, mode))
        p.close(); m.close()


def test_count_kernel_routes_wide_ranges_to_split(restore_tuning):
    """Integer columns whose value range exceeds the histogram limit go to the
    split path when they are read as rows and values; packed, their bins are
    the ~2,400 distinct values each holds, and the count kernel takes them.
    Results are the same either way."""
    say("count kernel routing")
    w = _random_workload(12, 6000, 10, 6, 0.4, lambda r, n: r.integers(-20000, 20000, n))
    for packed, counted in ((0, 0), (1, 10)):
        _capi.set_tuning(count_packed=packed)
        m, p = _plan_of(w, "mix")
        p.run(2)
        info = p.info()
        assert info["count_columns"] == counted and info["split_columns"] == 10 - counted
        assert_parity(p.results(with_stats=True), oracle_stats(w, "mix"))
        p.close(); m.close()


def test_count_kernel_routes_too_many_slices_to_split(restore_tuning):
    """300 targets x 1,000 values in 16 KiB: a workgroup holds 3 targets' rows,
    100 slices per column, more than 64: the columns take the split path."""
    say("count kernel too many slices")
    _capi.set_tuning(count_lds_bytes=16 * 1024, count_packed=0)
    w = _random_workload(14, 6000, 4, 300, 0.5, lambda r, n: r.integers(1, 1001, n))
    m, p = _plan_of(w, "min")
    p.run(2)
    info = p.info()
    assert info["count_columns"] == 0 and info["split_columns"] == 4
    assert_parity(p.results(with_stats=True), oracle_stats(w, "min"))
    p.close(); m.close()


@pytest.mark.parametrize("lds", [76 * 1024, 24 * 1024])
def test_count_kernel_target_slices(lds, restore_tuning):
    """200 targets x 400 values do not fit one workgroup: 3 slices of 67
    targets in 76 KiB, 9 slices of 23 in 24 KiB, each reading the column."""
    say(f"count kernel slices lds={lds}")
    _capi.set_tuning(count_lds_bytes=lds)
    w = _random_workload(15, 8000, 6, 200, 0.5, lambda r, n: r.integers(1, 401, n))
    for mode in ("min", "mix", "none"):
        m, p = _plan_of(w, mode)
        p.run(1)
        info = p.info()
        assert info["handed_over"] == 0 and info["count_columns"] == 6
        assert info["count_blocks"] == 6 * (3 if lds == 76 * 1024 else 9)
        assert_parity(p.results(with_stats=True), oracle_stats(w, mode))
        p.close(); m.close()


def test_count_kernel_extreme_values(restore_tuning):
    """Values at the ends of the 16-bit range of the column metadata:
    -32768 .. -32700 and 32700 .. 32767 in the count kernel; both ends in one
    column (65,536 values) in the split path when read as rows and values, in
    the count kernel when packed (its ~70 distinct values are the bins)."""
    say("count kernel extreme values")
    w = _random_workload(16, 5000, 3, 6, 0.4, lambda r, n: r.integers(0, 69, n))
    d, ip = w["data"].copy(), w["indptr"]
    d[ip[0]:ip[1]] = np.where(d[ip[0]:ip[1]] == 0, 0, d[ip[0]:ip[1]] - 32768)
    d[ip[0]] = -32768
    d[ip[1]:ip[2]] = np.where(d[ip[1]:ip[2]] == 0, 0, d[ip[1]:ip[2]] + 32699)
    d[ip[1]] = 32767
    d[ip[2]] = -32768
    d[ip[2] + 1] = 32767
    w = dict(w, data=d)
    for packed, counted in ((0, 2), (1, 3)):
        _capi.set_tuning(count_packed=packed)
        for mode in ("min", "mix", "none"):
            m, p = _plan_of(w, mode)
            p.run(1)
            info = p.info()
            assert info["count_columns"] == counted and info["split_columns"] == 3 - counted
            assert_parity(p.results(with_stats=True), oracle_stats(w, mode))
            p.close(); m.close()


def test_plan_without_stats_keeps_u1_and_p_only():
    say("plan without stats")
    w = _random_workload(13, 3000, 6, 4, 0.3, lambda r, n: r.integers(1, 20, n))
    m, p = _plan_of(w, "min", stats=False)
    p.run(1)
    U1, P = p.results()
    o = oracle_stats(w, "min")
    assert np.array_equal(U1, o[0]) and rel_close(P, o[1], rtol=P_RTOL)
    with pytest.raises(_capi.HpdexError, match="HPDEX_PLAN_STATS"):
        p.results(with_stats=True)
    st = p.stage_times(2)
    assert st["count"] > 0
    p.close(); m.close()


def test_plans_plain_batched_bit_identical():
    say("plans")
    w = _random_workload(8, 8000, 200, 12, 0.15, lambda r, n: r.negative_binomial(1, 0.3, n) + 1)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    oneshot = _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], 12, opts,
                            with_stats=True)
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    for form in (0, 1):
        p = _capi.Plan(m, w["group_id"], 12, opts, form=form, stats=True)
        p.run(3)
        res = p.results(with_stats=True)
        for a, b in zip(res, oneshot):
            assert np.array_equal(a, b)
        p.close()
    m.close()


def test_histogram_workload_full_size():
    """The 150,000 x 8,000 integer-count workload: batched == plain bit for
    bit, structural properties everywhere, and oracle parity on every one of
    the 2.4M (gene, target) pairs."""
    say("histogram workload generate")
    w = synthetic.histogram_case()
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    say("histogram workload batched")
    pb = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1, stats=True)
    pb.run(2)
    U1, P, tie, n1, n2 = pb.results(with_stats=True)
    say("histogram workload plain")
    pp = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=0, stats=True)
    pp.run(1)
    res_plain = pp.results(with_stats=True)
    for a, b in zip((U1, P, tie, n1, n2), res_plain):
        assert np.array_equal(a, b)
    assert np.all(U1 >= 0) and np.all(U1 <= n1 * n2)
    assert np.array_equal(2 * U1, np.round(2 * U1))
    assert np.all((P >= 0) & (P <= 1))
    assert np.all(n1 == 15000)
    sizes = np.bincount(w["group_id"][w["group_id"] > 0], minlength=w["n_targets"] + 1)[1:]
    assert np.array_equal(n2.reshape(-1, w["n_targets"])[0], sizes.astype(float))
    say("histogram workload oracle on every column")
    assert_parity((U1, P, tie, n1, n2), oracle_stats(w, "min"))
    pb.close(); pp.close(); m.close()


@pytest.fixture(scope="module")
def float_workload():
    say("float workload generate")
    return synthetic.float_case()


@pytest.mark.parametrize("mode", ["min", "mix"])
def test_float_workload_full_size(float_workload, mode):
    """The 400,000 x 18,000 log1p(CP10k) workload in zero_handling min and mix
    (every column on the sort path, split staged in LDS, target windows):
    batched == plain bit for bit, structural properties everywhere, oracle
    parity on a column sample that includes the densest genes."""
    w = float_workload
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=ZH[mode])
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    say(f"float workload batched {mode}")
    pb = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1, stats=True)
    pb.run(2)
    U1, P, tie, n1, n2 = pb.results(with_stats=True)
    assert pb.info()["split_columns"] == w["n_cols"]
    say("float workload plain")
    pp = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=0, stats=True)
    pp.run(1)
    res_plain = pp.results(with_stats=True)
    for a, b in zip((U1, P, tie, n1, n2), res_plain):
        assert np.array_equal(a, b)
    assert np.all(U1 >= 0) and np.all(U1 <= n1 * n2)
    assert np.array_equal(2 * U1, np.round(2 * U1))
    assert np.all((P >= 0) & (P <= 1))
    assert np.all(n1 == np.count_nonzero(w["group_id"] == 0))
    sizes = np.bincount(w["group_id"][w["group_id"] > 0], minlength=w["n_targets"] + 1)[1:]
    assert np.array_equal(n2.reshape(-1, w["n_targets"])[0], sizes.astype(float))
    nnz = np.diff(w["indptr"])
    cols = np.unique(np.concatenate([np.argsort(nnz)[-8:], np.arange(0, w["n_cols"], 997)]))
    sub = synthetic.subset_columns(w, cols)
    say(f"float workload oracle on {len(cols)} columns")
    o = oracle_stats(sub, mode)
    T = w["n_targets"]
    g = tuple(a.reshape(-1, T)[cols].reshape(-1) for a in (U1, P, tie, n1, n2))
    assert_parity(g, o)
    pb.close(); pp.close(); m.close()


def test_exact_workload_full_size():
    """The dense 150,000 x 6,000 workload with 2,000 targets of 3-25 cells and
    a 60-cell reference, exact p in FP64 (tie correction on, continuity off):
    batched == plain bit for bit, structural properties on every pair, and
    oracle parity (U1 and tie bit-exact, p within 1e-12 of the FP64 DP) on a
    column sample -- every column holds every target."""
    say("exact workload generate")
    w = synthetic.exact_case()
    opts = _capi.make_opts(tie_correction=True, use_continuity=False, zero_handling=1, method=1)
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    say("exact workload batched")
    pb = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1, stats=True)
    pb.run(2)
    U1, P, tie, n1, n2 = pb.results(with_stats=True)
    say("exact workload plain")
    pp = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=0, stats=True)
    pp.run(1)
    for a, b in zip((U1, P, tie, n1, n2), pp.results(with_stats=True)):
        assert np.array_equal(a, b)
    pb.close(); pp.close(); m.close()
    T = w["n_targets"]
    sizes = np.bincount(w["group_id"][w["group_id"] > 0], minlength=T + 1)[1:]
    assert np.all(n1 == 60)
    assert np.array_equal(n2.reshape(-1, T), np.broadcast_to(sizes.astype(float), (w["n_cols"], T)))
    assert np.all(U1 >= 0) and np.all(U1 <= n1 * n2)
    assert np.array_equal(2 * U1, np.round(2 * U1))
    assert np.all((P > 0) & (P <= 1))
    # float32 normals are nearly tie-free: few pairs carry a tie term, each
    # a sum of t^3 - t over runs of t >= 2 (at least 6)
    assert np.all(tie >= 0) and np.array_equal(tie, np.round(tie))
    assert np.all((tie == 0) | (tie >= 6)) and np.mean(tie > 0) < 0.05
    cols = np.arange(0, w["n_cols"], 40)
    sub = synthetic.subset_columns(w, cols)
    say(f"exact workload oracle on {len(cols)} columns")
    o = oracle_stats(sub, "min", method="exact", cc=False)
    g = tuple(a.reshape(-1, T)[cols].reshape(-1) for a in (U1, P, tie, n1, n2))
    assert_parity(g, o)


def test_plan_kernel_ms_keeps_results():
    """Timing each kernel alone (back-to-back launches between HIP events)
    leaves the plan's results and state those of a normal run, on a job with
    count-path and sort-path columns."""
    say("plan kernel_ms")
    rng = np.random.default_rng(29)
    R, T = 6000, 12
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.6, 0.3, 0.5, 0.2]))
    data = rng.integers(1, 20, len(rows)).astype(np.float32)
    data[indptr[2]:] += np.float32(0.5)   # two float columns: the sort path
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:1500]] = 0
    gid[perm[1500:]] = rng.integers(1, T + 1, R - 1500)
    w = dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)
    m, p = _plan_of(w, "min", form=1)
    p.run(1)
    want = p.results(with_stats=True)
    for k in ("count", "split", "reference", "target"):
        assert p.kernel_ms(k, iters=5) > 0.0
        for a, b in zip(p.results(with_stats=True), want):
            assert np.array_equal(a, b)
    p.run(1)
    for a, b in zip(p.results(with_stats=True), want):
        assert np.array_equal(a, b)
    assert_parity(want, oracle_stats(w, "min"))
    with pytest.raises(ValueError):
        p.kernel_ms("nothing")
    p.close()
    m.close()


@pytest.mark.parametrize("mode", ["min", "max", "mix", "none"])
def test_float_targets_of_65_to_192_keys(mode):
    """Float targets of 65 to 192 keys (ranked by the window kernel without a
    sort, tie positions from the wave's table), beside targets of more keys
    (the sorted pass) and of at most 64 (windows); many equal keys inside a
    target.  U1, tie and sizes bit-exact, p within 1e-12 of the oracle."""
    say(f"targets of 65-192 keys, {mode}")
    rng = np.random.default_rng(61)
    R = 20000
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.9, 0.6, 0.35, 0.2, 0.15]))
    data = np.round(rng.normal(0, 1.5, len(rows)), 1).astype(np.float32)
    sizes = [90, 120, 140, 180, 260, 330, 60, 75, 400, 700, 2000]
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:3000]] = 0
    k = 3000
    for t, n in enumerate(sizes, start=1):
        gid[perm[k:k + n]] = t
        k += n
    T = len(sizes)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True,
                           zero_handling={"none": 0, "min": 1, "max": 2, "mix": 3}[mode])
    g = _capi.mwu_csc(data, rows, indptr, gid, T, opts, n_rows=R, with_stats=True)
    o = oracle.mwu(data, rows, indptr, gid, T, n_rows=R, tie_correction=True,
                   use_continuity=True, zero_handling=mode, with_stats=True)
    for a, b in zip((g[0], g[2], g[3], g[4]), (o[0], o[2], o[3], o[4])):
        assert np.array_equal(a, b)
    assert rel_close(g[1], o[1], rtol=1e-12)
