"""Device ingest (hpdex_matrix_create_from): CSR, dense and 32-bit-index CSC
input become on the device exactly the CSC that scipy builds on the host
(tocsc / csc_matrix, src/hpdex/de.py:103-109), and every result computed on
them equals the result on the host-converted CSC bit for bit."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same_csc(dm, ref):
    """The resident matrix equals scipy's CSC `ref` (values as float64)."""
    ref = ref.tocsc()
    data, rows, indptr = dm.download()
    assert dm.n_rows == ref.shape[0] and dm.n_cols == ref.shape[1]
    assert np.array_equal(indptr, ref.indptr.astype(np.int64))
    assert np.array_equal(rows, ref.indices.astype(np.int32))
    assert np.array_equal(data, ref.data.astype(np.float64), equal_nan=True)


def _matrix(seed, R, C, dens, dtype, zeros=True, nan=False):
    rng = np.random.default_rng(seed)
    X = sp.random(R, C, density=dens, format="csr", random_state=seed, dtype=np.float64)
    X.data = np.floor(X.data * 50) - (5 if zeros else -1)      # explicit zeros and negatives
    if nan:
        X.data[rng.random(len(X.data)) < 0.01] = np.nan
    X.data = X.data.astype(dtype)
    X.sort_indices()
    return X


@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.float64, np.uint8])
@pytest.mark.parametrize("bits", [32, 64])
def test_csr_to_csc_equals_scipy(dtype, bits):
    say(f"csr -> csc {np.dtype(dtype).name} int{bits}")
    nan = np.issubdtype(dtype, np.floating)
    X = _matrix(41, 3000, 170, 0.07, dtype, zeros=not np.issubdtype(dtype, np.unsignedinteger),
                nan=nan)
    X.indices = X.indices.astype(np.int64 if bits == 64 else np.int32)
    X.indptr = X.indptr.astype(np.int64 if bits == 64 else np.int32)
    dm = _capi.DeviceMatrix.from_scipy(X)
    assert dm.nnz == X.nnz
    _same_csc(dm, X)
    dm.close()


def test_csr_edge_shapes():
    """Empty rows and columns, a single row, an empty matrix."""
    say("csr edge shapes")
    X = sp.csr_matrix(np.array([[0, 0, 3, 0], [0, 0, 0, 0], [1, 0, 2, 0]], dtype=np.float32))
    for m in (X, sp.csr_matrix((5, 7), dtype=np.float32), sp.csr_matrix(np.ones((1, 9), np.float32))):
        dm = _capi.DeviceMatrix.from_scipy(m)
        _same_csc(dm, m)
        dm.close()


def test_csr_bad_column_index_is_an_error():
    say("csr bad column")
    X = _matrix(42, 200, 30, 0.1, np.float32)
    X.indices[5] = 30          # one past the last column
    with pytest.raises(_capi.HpdexError, match="column index out of range"):
        _capi.DeviceMatrix.from_scipy(X)


@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.float64])
def test_dense_to_csc_equals_scipy(dtype):
    """Zeros (and -0.0) dropped, NaN kept, as csc_matrix(X) does."""
    say(f"dense -> csc {np.dtype(dtype).name}")
    X = _matrix(43, 2100, 130, 0.2, dtype, nan=np.issubdtype(dtype, np.floating)).toarray()
    if np.issubdtype(dtype, np.floating):
        X[3, :10] = -0.0
    dm = _capi.DeviceMatrix.from_dense(X)
    _same_csc(dm, sp.csc_matrix(X))
    dm.close()


def test_csc_int32_indices_and_rows_out_of_range():
    """32-bit CSC uploads directly; rows outside [0, R) read as -1, as the
    64-bit upload marks them."""
    say("csc int32")
    X = _matrix(44, 1000, 60, 0.1, np.float32).tocsc()
    X.indices = X.indices.astype(np.int32)
    X.indices[7] = 5000
    X.indices[8] = -3
    dm32 = _capi.DeviceMatrix.from_scipy(X)
    dm64 = _capi.DeviceMatrix(X.data, X.indices.astype(np.int64), X.indptr.astype(np.int64), 1000)
    a, b = dm32.download(), _capi.DeviceMatrix.download(dm64)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert a[1][7] == -1 and a[1][8] == -1
    dm32.close(); dm64.close()


def _plan_results(dm, gid, T, mode):
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=mode)
    p = _capi.Plan(dm, gid, T, opts, form=1, stats=True)
    p.run(1)
    r = p.results(with_stats=True)
    p.close()
    return r


@pytest.mark.parametrize("layout", ["csr", "dense"])
def test_results_identical_across_ingest_routes(layout):
    """Count and sort path (integer and float columns), all zero modes: the
    matrix built on the device gives the bit-identical U1, tie, sizes and p of
    the host-converted CSC."""
    say(f"results across ingest {layout}")
    rng = np.random.default_rng(45)
    R, C, T = 4000, 48, 7
    X = _matrix(45, R, C, 0.15, np.float32)
    X.data[X.indices >= 24] += 0.25                # half the genes non-integer: sort path
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    csc = X.tocsc()
    ref = _capi.DeviceMatrix(csc.data, csc.indices.astype(np.int64), csc.indptr.astype(np.int64), R)
    dev = _capi.DeviceMatrix.from_scipy(X) if layout == "csr" else \
        _capi.DeviceMatrix.from_dense(X.toarray())
    if layout == "dense":       # csc_matrix(X) drops explicit zeros: compare with that
        ref.close()
        dcsc = sp.csc_matrix(X.toarray())
        ref = _capi.DeviceMatrix(dcsc.data, dcsc.indices.astype(np.int64),
                                 dcsc.indptr.astype(np.int64), R)
    for mode in (0, 1, 2, 3):
        a = _plan_results(dev, gid, T, mode)
        b = _plan_results(ref, gid, T, mode)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), mode
    dev.close(); ref.close()


def test_de_matrix_equals_de_csc():
    say("de_matrix")
    rng = np.random.default_rng(46)
    X = _matrix(46, 3000, 40, 0.2, np.float32, zeros=False)
    gid = rng.integers(-1, 5, 3000).astype(np.int32)
    csc = X.tocsc()
    a = _capi.de_csc(csc.data, csc.indices.astype(np.int64), csc.indptr.astype(np.int64), gid, 4,
                     n_rows=3000)
    dm = _capi.DeviceMatrix.from_scipy(X)
    b = _capi.de_matrix(dm, gid, 4)
    dm.close()
    assert all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


def test_histogram_workload_as_csr():
    """The 150,000 x 8,000 workload handed over as CSR with int32 indices
    (AnnData's layout) becomes its CSC on the device."""
    say("150k x 8000 as csr")
    w = synthetic.histogram_case()
    csc = sp.csc_matrix((w["data"], w["indices"], w["indptr"]), shape=(w["n_rows"], w["n_cols"]))
    csr = csc.tocsr()
    csr.indices = csr.indices.astype(np.int32)
    csr.indptr = csr.indptr.astype(np.int32)
    dm = _capi.DeviceMatrix.from_scipy(csr)
    data, rows, indptr = dm.download()
    dm.close()
    assert np.array_equal(indptr, w["indptr"])
    assert np.array_equal(rows, w["indices"].astype(np.int32))
    assert np.array_equal(data, w["data"].astype(np.float64))
