"""hpdex_mwu_csc_noalloc through the C-ABI: the same results as
hpdex_mwu_csc, bit for bit, on two calls in a row, the second with a smaller
input (the scratch keeps the first call's contents), on the count path and
the sort path; a job above the scratch is refused before anything runs."""
import sys

import numpy as np
import pytest

from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _float_case(seed, R, C, T):
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, 0.3))
    data = np.round(rng.normal(0, 2, len(rows)), 2).astype(np.float32)
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    return dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)


def _both(w, opts):
    a = _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"], opts,
                      n_rows=w["n_rows"])
    b = _capi.mwu_csc_noalloc(w["data"], w["indices"], w["indptr"], w["group_id"],
                              w["n_targets"], opts, n_rows=w["n_rows"])
    return a, b


def test_noalloc_equals_mwu_csc_twice_second_smaller():
    big = synthetic.small_case()                              # 2,000 x 500, count path
    small = synthetic.subset_columns(big, np.arange(0, 500, 7))
    small["group_id"] = np.where(big["group_id"] > 2, -1, big["group_id"]).astype(np.int32)
    small["n_targets"] = 2
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    for name, w in (("2000x500", big), ("2000x72", small)):
        say(f"noalloc count path {name}")
        (U1a, Pa), (U1b, Pb) = _both(w, opts)
        assert np.array_equal(U1a, U1b) and np.array_equal(Pa, Pb)
    opts = _capi.make_opts(tie_correction=True, use_continuity=False, zero_handling=3)
    for name, w in (("3000x200", _float_case(3, 3000, 200, 6)),
                    ("1500x40", _float_case(4, 1500, 40, 3))):
        say(f"noalloc sort path {name}")
        (U1a, Pa), (U1b, Pb) = _both(w, opts)
        assert np.array_equal(U1a, U1b) and np.array_equal(Pa, Pb)


def test_noalloc_exact_and_errors():
    w = synthetic.exact_case(n_cells=3000, n_genes=30, n_ref=60, n_targets=40)
    for mode in (1, 0):
        say(f"noalloc exact zero_handling={mode}")
        opts = _capi.make_opts(tie_correction=True, use_continuity=False, zero_handling=mode,
                               method=1)
        (U1a, Pa), (U1b, Pb) = _both(w, opts)
        assert np.array_equal(U1a, U1b) and np.array_equal(Pa, Pb)
    say("noalloc errors")
    # the reference's messages, as hpdex_mwu_csc gives them
    gid = w["group_id"].copy()
    gid[gid == 0] = -1
    gid[np.argmax(w["group_id"] >= 0)] = 0
    opts = _capi.make_opts(zero_handling=1)
    with pytest.raises(_capi.HpdexError, match="Sample too small for reference at column 0"):
        _capi.mwu_csc_noalloc(w["data"], w["indices"], w["indptr"], gid, 40, opts,
                              n_rows=w["n_rows"])
    with pytest.raises(_capi.HpdexError, match="n_targets must be >= 1"):
        _capi.mwu_csc_noalloc(w["data"], w["indices"], w["indptr"], gid, 0, opts)
    # 3,000 columns take about 100 MiB of per-column tables: refused
    many = _float_case(5, 500, 3000, 2)
    with pytest.raises(_capi.HpdexError, match="more than its scratch of 67108864 bytes"):
        _capi.mwu_csc_noalloc(many["data"], many["indices"], many["indptr"], many["group_id"], 2,
                              opts, n_rows=500)
    # and the scratch still serves the next call
    small = _float_case(6, 800, 20, 3)
    (U1a, Pa), (U1b, Pb) = _both(small, opts)
    assert np.array_equal(U1a, U1b) and np.array_equal(Pa, Pb)
