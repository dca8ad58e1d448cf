"""Compact uploads: a float32 CSC matrix of small whole-number values (UMI
counts) crosses PCIe in 4 bytes per entry -- the row's low 16 bits and the
value as a 16-bit integer -- and is decoded on the device into the same
resident matrix (rows and values bit for bit) as the 8-byte upload.  A column
holding an entry that does not qualify (another value, -0.0, NaN, a row out of
range, out of order, repeated or 2^16 above the one before) is written as it
is after the decode, wherever in the copy that entry comes; a matrix with more
than an eighth of its entries in such columns is uploaded as it is."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from fixtures import rel_close
from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def counts(seed, R, dens, top=60):
    """Integer counts as float32 in CSC (rows ascending), int64 indices."""
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.asarray(dens, float))
    data = (rng.integers(1, top, len(rows))).astype(np.float32)
    return data, rows.astype(np.int64), indptr.astype(np.int64)


def expect_rows(rows, R):
    r = np.asarray(rows, dtype=np.int64)
    return np.where((r >= 0) & (r < R), r, -1).astype(np.int32)


def check_resident(m, data, rows, indptr, R, compact, patched=0):
    assert m.compact_upload() == (compact, patched)
    d, r, p = m.download()
    assert np.array_equal(p, indptr)
    assert np.array_equal(r, expect_rows(rows, R))
    # values as the device holds them (float32), compared bit for bit
    assert np.array_equal(d.astype(np.float32).view(np.uint32),
                          np.asarray(data, np.float32).view(np.uint32))


@pytest.mark.parametrize("R", [40_000, 300_000], ids=["rows-16bit", "rows-wrap"])
def test_compact_upload_round_trip(R):
    """Columns from empty to dense (many wraps of the low row bits when R >
    2^16, columns longer than one 1,024-entry decode chunk), values up to
    65,535 and explicit zeros: the resident matrix equals the source."""
    say(f"compact round trip R={R}")
    dens = [0.0, 1.0, 0.5, 0.2, 0.05, 0.0, 0.9] + [0.3] * (4_000_000 // R)
    data, rows, indptr = counts(3, R, dens)
    data[::97] = 0.0
    data[5::101] = 65535.0
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    check_resident(m, data, rows, indptr, R, True)
    m.close()


def test_compact_upload_int32_source_and_results():
    """A scipy CSC with int32 indices through hpdex_matrix_create_from: the
    compact form, and Mann-Whitney results equal to the oracle's."""
    say("compact int32 source")
    w = synthetic.histogram_case(n_cells=60_000, n_genes=300)
    X = sp.csc_matrix((w["data"], w["indices"].astype(np.int32), w["indptr"].astype(np.int32)),
                      shape=(w["n_rows"], w["n_cols"]))
    m = _capi.DeviceMatrix.from_scipy(X)
    assert m.compact_upload()[0]
    check_resident(m, w["data"], w["indices"], w["indptr"], w["n_rows"], True,
                   m.compact_upload()[1])
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    p = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1)
    p.run(1)
    U1, P = p.results()
    oU1, oP = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                         n_rows=w["n_rows"], tie_correction=True, use_continuity=True,
                         zero_handling="min", threads=8)
    assert np.array_equal(U1, oU1) and rel_close(P, oP, rtol=1e-12)
    p.close()
    m.close()


def _spoil(kind, data, rows, indptr, at):
    data, rows = data.copy(), rows.copy()
    c = int(np.searchsorted(indptr, at, side="right") - 1)
    e = max(at, int(indptr[c]) + 1)   # an entry with a predecessor in its column
    if kind == "fraction":
        data[at] = 2.5
    elif kind == "too-large":
        data[at] = 65536.0
    elif kind == "negative":
        data[at] = -3.0
    elif kind == "minus-zero":
        data[at] = -0.0
    elif kind == "nan":
        data[at] = np.nan
    elif kind == "row-out-of-range":
        rows[at] = -1
    elif kind == "row-repeated":
        rows[e] = rows[e - 1]
    elif kind == "row-descending":
        rows[e - 1], rows[e] = rows[e], rows[e - 1]
    return data, rows


@pytest.mark.parametrize("kind", ["fraction", "too-large", "negative", "minus-zero", "nan",
                                  "row-out-of-range", "row-repeated", "row-descending"])
@pytest.mark.parametrize("where", ["first", "late"])
def test_columns_that_do_not_qualify_are_written_as_they_are(kind, where):
    """One entry that does not qualify, in the first staging chunk or in the
    last (after chunks were sent compactly): its column is written as it is,
    the others decoded, and the resident matrix equals the source."""
    say(f"compact patch {kind} {where}")
    R = 120_000
    data, rows, indptr = counts(4, R, [0.6] * 160)   # ~11.5M entries: three 16 MB chunks
    nnz = len(data)
    at = 5 if where == "first" else nnz - 1000
    data, rows = _spoil(kind, data, rows, indptr, at)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    assert m.compact_upload() == (True, 1)
    d, r, p = m.download()
    assert np.array_equal(r, expect_rows(rows, R)) and np.array_equal(p, indptr)
    assert np.array_equal(d.astype(np.float32).view(np.uint32),
                          data.astype(np.float32).view(np.uint32))
    m.close()


def test_float_values_upload_as_they_are():
    """Values that are not whole numbers everywhere: more than an eighth of
    the entries in columns that do not qualify, so the compact copy stops and
    the matrix is uploaded as it is."""
    say("compact float values")
    R = 120_000
    data, rows, indptr = counts(7, R, [0.6] * 40)
    data = data + np.float32(0.25)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    check_resident(m, data, rows, indptr, R, False)
    m.close()


def test_row_gap_of_two_to_the_sixteen_is_patched():
    """Rows 2^16 apart in one column cannot be told from a gap of 0 by their
    low bits: that column is written as it is."""
    say("compact gap 2^16")
    R = 200_000
    data, rows, indptr = counts(5, R, [0.4] * 30)
    c0 = slice(indptr[3], indptr[4])
    col = np.array([10, 10 + 65536, 10 + 2 * 65536], dtype=np.int64)
    rows = np.concatenate([rows[:indptr[3]], col, rows[indptr[4]:]])
    data = np.concatenate([data[:indptr[3]], np.ones(3, np.float32), data[indptr[4]:]])
    indptr = indptr.copy()
    indptr[4:] += 3 - (c0.stop - c0.start)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    check_resident(m, data, rows, indptr, R, True, 1)
    m.close()


def test_small_and_plain_copies_upload_as_they_are():
    """Copies below the staging threshold (4 MB) go as they are."""
    say("compact small")
    data, rows, indptr = counts(6, 5000, [0.5] * 20)
    m = _capi.DeviceMatrix(data, rows, indptr, 5000)
    check_resident(m, data, rows, indptr, 5000, False)
    m.close()


def test_one_shot_and_session_jobs_on_compact_uploads():
    """The one-shot call and a session job that brings its int32 CSC take the
    compact form on the way; results equal the oracle's."""
    say("compact one-shot and session")
    w = synthetic.histogram_case(n_cells=60_000, n_genes=400)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    oU1, oP = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                         n_rows=w["n_rows"], tie_correction=True, use_continuity=True,
                         zero_handling="min", threads=8)
    U1, P = _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                          opts, n_rows=w["n_rows"])
    assert np.array_equal(U1, oU1) and rel_close(P, oP, rtol=1e-12)
    s = _capi.Session(8 << 30, slots=2)
    t = s.submit(w["group_id"], w["n_targets"], opts, data=w["data"],
                 indices=w["indices"].astype(np.int32), indptr=w["indptr"].astype(np.int32),
                 n_rows=w["n_rows"])
    U1s, Ps, _ = s.wait(t)
    s.close()
    assert np.array_equal(U1s, oU1) and rel_close(Ps, oP, rtol=1e-12)
