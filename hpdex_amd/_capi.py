"""ctypes view of the C-ABI in include/hpdex_mi355x.h.

Used by the plan/serving layer, the tests and bench.py to call the native
library directly.  Loading fails loudly if the library has not been built:
there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhpdex_mi355x.so")

DTYPE_CODES = {
    np.dtype(np.int8): 0, np.dtype(np.uint8): 1, np.dtype(np.int16): 2,
    np.dtype(np.uint16): 3, np.dtype(np.int32): 4, np.dtype(np.uint32): 5,
    np.dtype(np.int64): 6, np.dtype(np.uint64): 7, np.dtype(np.float32): 8,
    np.dtype(np.float64): 9,
}

# Symbols declared in include/hpdex_mi355x.h (checked by the CPU tests).
EXPORTED = (
    "hpdex_mwu_csc", "hpdex_mwu_csc_stats", "hpdex_group_mean_csc",
    "hpdex_matrix_create", "hpdex_matrix_destroy", "hpdex_matrix_device_bytes",
    "hpdex_plan_create", "hpdex_plan_destroy", "hpdex_plan_run", "hpdex_plan_results",
    "hpdex_get_tuning", "hpdex_set_tuning", "hpdex_version", "hpdex_device_info",
    "hpdex_bandwidth_probe", "hpdex_plan_get_info", "hpdex_plan_stage_times",
    "hpdex_session_create", "hpdex_session_submit", "hpdex_session_wait", "hpdex_session_peak",
    "hpdex_session_destroy", "hpdex_fdr_bh", "hpdex_de_csc", "hpdex_matrix_create_from",
    "hpdex_matrix_shape", "hpdex_matrix_download", "hpdex_de_matrix",
    "hpdex_de_csc_device_bytes", "hpdex_mwu_csc_noalloc",
    "hpdex_batch_create", "hpdex_batch_run", "hpdex_batch_get_info", "hpdex_batch_destroy",
    "hpdex_session_batches", "hpdex_pden_matrix", "hpdex_matrix_ref_cache_info",
    "hpdex_matrix_ref_cache_clear", "hpdex_matrix_upload_form", "hpdex_plan_kernel_ms",
)

LAYOUT_CSC, LAYOUT_CSR, LAYOUT_DENSE = 0, 1, 2


class MatrixSrc(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_int32), ("dtype", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("indices", ctypes.c_void_p), ("indptr", ctypes.c_void_p),
                ("index_bits", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("n_rows", ctypes.c_size_t), ("n_cols", ctypes.c_size_t), ("nnz", ctypes.c_size_t)]


class MwuOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "ref_sorted", "tar_sorted", "tie_correction", "use_continuity", "fast_norm",
        "zero_handling", "alternative", "method")]


class PdenOpts(ctypes.Structure):
    _fields_ = [("tie_correction", ctypes.c_int32), ("continuity_correction", ctypes.c_int32),
                ("use_asymptotic", ctypes.c_int32), ("use_hist", ctypes.c_int32),
                ("max_bins", ctypes.c_int64), ("chunk_genes", ctypes.c_int64),
                ("clip_value", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("hist_max_bins", "size_ordered", "count_lds_bytes",
                                              "count_block", "count_packed", "split_arrival",
                                              "ref_lds_keys", "ref_cache")]


class PlanInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "count_columns", "count_nnz", "split_columns", "handed_over", "count_blocks",
        "count_packed_columns", "ref_cached_columns")]


class SessionOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("slots", ctypes.c_int32)]


class Job(ctypes.Structure):
    _fields_ = [("matrix", ctypes.c_void_p), ("dtype", ctypes.c_int32),
                ("data", ctypes.c_void_p), ("indices", ctypes.c_void_p),
                ("indptr", ctypes.c_void_p), ("n_rows", ctypes.c_size_t),
                ("n_cols", ctypes.c_size_t), ("nnz", ctypes.c_size_t),
                ("group_id", ctypes.c_void_p), ("n_targets", ctypes.c_size_t),
                ("opts", MwuOpts), ("U1_out", ctypes.c_void_p), ("P_out", ctypes.c_void_p),
                ("index_bits", ctypes.c_int32), ("layout", ctypes.c_int32)]


class JobTimes(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("footprint_bytes", ctypes.c_size_t),
                ("upload_ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("kernels_ms", ctypes.c_double), ("finish_ms", ctypes.c_double)]


class HpdexError(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m hpdex_amd.build` "
                "(the MI355X backend has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, ci, cp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_char_p
        L.hpdex_mwu_csc.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz, ctypes.POINTER(MwuOpts),
                                    ci, vp, vp, cp, sz]
        L.hpdex_mwu_csc_stats.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz,
                                          ctypes.POINTER(MwuOpts), ci, vp, vp, vp, vp, vp, cp, sz]
        L.hpdex_group_mean_csc.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz, ci, ci, ci, vp,
                                           cp, sz]
        L.hpdex_matrix_create.argtypes = [ci, vp, vp, vp, sz, sz, sz, ctypes.POINTER(vp), cp, sz]
        L.hpdex_matrix_destroy.argtypes = [vp]
        L.hpdex_matrix_destroy.restype = None
        L.hpdex_matrix_device_bytes.argtypes = [vp]
        L.hpdex_matrix_device_bytes.restype = sz
        L.hpdex_matrix_ref_cache_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                                  ctypes.POINTER(sz)]
        L.hpdex_matrix_ref_cache_clear.argtypes = [vp]
        L.hpdex_matrix_ref_cache_clear.restype = None
        L.hpdex_matrix_upload_form.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.hpdex_matrix_upload_form.restype = ctypes.c_int
        L.hpdex_plan_create.argtypes = [vp, vp, sz, ctypes.POINTER(MwuOpts), ci,
                                        ctypes.POINTER(vp), cp, sz]
        L.hpdex_plan_destroy.argtypes = [vp]
        L.hpdex_plan_destroy.restype = None
        L.hpdex_plan_run.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), cp, sz]
        L.hpdex_plan_results.argtypes = [vp, vp, vp, vp, vp, vp, cp, sz]
        L.hpdex_get_tuning.argtypes = [ctypes.POINTER(Tuning)]
        L.hpdex_set_tuning.argtypes = [ctypes.POINTER(Tuning), cp, sz]
        L.hpdex_version.restype = cp
        L.hpdex_device_info.argtypes = [ctypes.POINTER(ci), ctypes.POINTER(ci), cp, sz]
        L.hpdex_bandwidth_probe.argtypes = [sz, ci, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.hpdex_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo), cp, sz]
        L.hpdex_session_create.argtypes = [ctypes.POINTER(SessionOpts), ctypes.POINTER(vp), cp, sz]
        L.hpdex_session_submit.argtypes = [vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_int64),
                                           cp, sz]
        L.hpdex_session_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(JobTimes), cp, sz]
        L.hpdex_session_peak.argtypes = [vp, ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_size_t)]
        L.hpdex_fdr_bh.argtypes = [vp, sz, vp, cp, sz]
        L.hpdex_de_csc.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz, ci, ci, ctypes.c_double, ci,
                                   vp, vp, vp, vp, vp, cp, sz]
        L.hpdex_session_destroy.argtypes = [vp]
        L.hpdex_session_destroy.restype = None
        L.hpdex_plan_stage_times.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.hpdex_plan_kernel_ms.argtypes = [vp, ci, ci, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.hpdex_plan_kernel_ms.restype = ci
        L.hpdex_matrix_create_from.argtypes = [ctypes.POINTER(MatrixSrc), ctypes.POINTER(vp), cp, sz]
        L.hpdex_matrix_shape.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz),
                                         ctypes.POINTER(sz)]
        L.hpdex_matrix_download.argtypes = [vp, vp, vp, vp, cp, sz]
        L.hpdex_de_matrix.argtypes = [vp, vp, sz, ci, ci, ctypes.c_double, vp, vp, vp, vp, vp, cp, sz]
        L.hpdex_pden_matrix.argtypes = [vp, vp, sz, ctypes.POINTER(PdenOpts), vp, vp, vp, vp, vp,
                                        cp, sz]
        L.hpdex_pden_matrix.restype = ci
        L.hpdex_mwu_csc_noalloc.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz,
                                            ctypes.POINTER(MwuOpts), ci, vp, vp, cp, sz]
        L.hpdex_mwu_csc_noalloc.restype = ci
        L.hpdex_batch_create.argtypes = [ctypes.POINTER(vp), sz, ctypes.POINTER(vp), cp, sz]
        L.hpdex_batch_create.restype = ci
        L.hpdex_batch_run.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_double), cp, sz]
        L.hpdex_batch_run.restype = ci
        L.hpdex_batch_get_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]
        L.hpdex_batch_get_info.restype = ci
        L.hpdex_session_batches.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int64)] * 3
        L.hpdex_session_batches.restype = ci
        L.hpdex_batch_destroy.argtypes = [vp]
        L.hpdex_batch_destroy.restype = None
        L.hpdex_de_csc_device_bytes.argtypes = [ci, sz, sz, sz, sz]
        L.hpdex_de_csc_device_bytes.restype = sz
        for name in ("hpdex_mwu_csc", "hpdex_mwu_csc_stats", "hpdex_group_mean_csc",
                     "hpdex_matrix_create", "hpdex_plan_create", "hpdex_plan_run",
                     "hpdex_plan_results", "hpdex_get_tuning", "hpdex_set_tuning",
                     "hpdex_device_info", "hpdex_bandwidth_probe", "hpdex_plan_get_info",
                     "hpdex_plan_stage_times", "hpdex_session_create", "hpdex_session_submit",
                     "hpdex_session_wait", "hpdex_session_peak", "hpdex_fdr_bh",
                     "hpdex_de_csc"):
            getattr(L, name).restype = ci
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def make_opts(tie_correction=True, use_continuity=False, zero_handling=3, alternative=2,
              method=2, ref_sorted=False, tar_sorted=False, fast_norm=False) -> MwuOpts:
    return MwuOpts(int(ref_sorted), int(tar_sorted), int(tie_correction), int(use_continuity),
                   int(fast_norm), int(zero_handling), int(alternative), int(method))


def _csc_args(data, indices, indptr):
    data = np.ascontiguousarray(data)
    if data.dtype not in DTYPE_CODES:
        raise HpdexError("Unsupported dtype for data")
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    return data, indices, indptr


def mwu_csc(data, indices, indptr, group_id, n_targets, opts: MwuOpts, n_rows=None,
            with_stats=False, threads=-1):
    """Call hpdex_mwu_csc(_stats) on host arrays.  Returns (U1, P) or
    (U1, P, tie, n1, n2)."""
    data, indices, indptr = _csc_args(data, indices, indptr)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    R = len(group_id) if n_rows is None else int(n_rows)
    C = len(indptr) - 1
    np_ = C * n_targets if n_targets > 0 else 0
    outs = [np.empty(max(np_, 1)) for _ in range(5 if with_stats else 2)]
    err = ctypes.create_string_buffer(1024)
    if with_stats:
        rc = lib().hpdex_mwu_csc_stats(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr),
                                       R, C, len(data), _p(group_id), int(n_targets),
                                       ctypes.byref(opts), threads, *[_p(o) for o in outs],
                                       err, len(err))
    else:
        rc = lib().hpdex_mwu_csc(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr), R, C,
                                 len(data), _p(group_id), int(n_targets), ctypes.byref(opts),
                                 threads, _p(outs[0]), _p(outs[1]), err, len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return tuple(o[:np_] for o in outs)


def mwu_csc_noalloc(data, indices, indptr, group_id, n_targets, opts: MwuOpts, n_rows=None,
                    threads=-1):
    """Call hpdex_mwu_csc_noalloc (no device allocation: the library's fixed
    scratch) on host arrays.  Returns (U1, P)."""
    data, indices, indptr = _csc_args(data, indices, indptr)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    R = len(group_id) if n_rows is None else int(n_rows)
    C = len(indptr) - 1
    np_ = C * n_targets if n_targets > 0 else 0
    outs = [np.empty(max(np_, 1)) for _ in range(2)]
    err = ctypes.create_string_buffer(1024)
    rc = lib().hpdex_mwu_csc_noalloc(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr),
                                     R, C, len(data), _p(group_id), int(n_targets),
                                     ctypes.byref(opts), threads, _p(outs[0]), _p(outs[1]), err,
                                     len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return tuple(o[:np_] for o in outs)


def group_mean_csc(data, indices, indptr, group_id, n_groups, include_zeros=True,
                   use_kahan=False, threads=-1):
    data, indices, indptr = _csc_args(data, indices, indptr)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    C = len(indptr) - 1
    out = np.empty(max(C * n_groups, 1))
    err = ctypes.create_string_buffer(1024)
    rc = lib().hpdex_group_mean_csc(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr),
                                    len(group_id), C, len(data), _p(group_id), int(n_groups),
                                    int(include_zeros), int(use_kahan), threads, _p(out),
                                    err, len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return out[:C * n_groups]


def fdr_bh(p):
    """Benjamini-Hochberg FDR on the device (hpdex_fdr_bh)."""
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
    out = np.empty(max(len(p), 1))
    err = ctypes.create_string_buffer(1024)
    if lib().hpdex_fdr_bh(_p(p), len(p), _p(out), err, len(err)) != 0:
        raise HpdexError(err.value.decode())
    return out[:len(p)]


def de_csc(data, indices, indptr, group_id, n_targets, n_rows=None, tie_correction=True,
           use_continuity=True, clip_value=20.0, threads=-1):
    """hpdex_de_csc: (U1, P, fold_change, log2_fold_change, fdr); fold changes
    target-major (t*C + c), the rest gene-major (c*T + t)."""
    data, indices, indptr = _csc_args(data, indices, indptr)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    R = len(group_id) if n_rows is None else int(n_rows)
    C = len(indptr) - 1
    n = C * int(n_targets)
    outs = [np.empty(max(n, 1)) for _ in range(5)]
    err = ctypes.create_string_buffer(1024)
    rc = lib().hpdex_de_csc(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr), R, C,
                            len(data), _p(group_id), int(n_targets), int(tie_correction),
                            int(use_continuity), float(clip_value), int(threads),
                            *[_p(o) for o in outs], err, len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return tuple(o[:n] for o in outs)


def de_csc_device_bytes(dtype, n_rows, n_cols, nnz, n_targets) -> int:
    """Upper bound of the device bytes one de_csc call on a matrix of this
    shape allocates (hpdex_de_csc_device_bytes; computed on the host)."""
    return int(lib().hpdex_de_csc_device_bytes(DTYPE_CODES[np.dtype(dtype)], int(n_rows),
                                                int(n_cols), int(nnz), int(n_targets)))


def _index_array(a):
    """Index array as given when it is int32 or int64 (no copy), else int64."""
    a = np.asarray(a)
    if a.dtype not in (np.dtype(np.int32), np.dtype(np.int64)):
        a = a.astype(np.int64)
    return np.ascontiguousarray(a)


def _data_array(a):
    a = np.asarray(a)
    if a.dtype not in DTYPE_CODES:
        raise HpdexError("Unsupported dtype for data")
    return np.ascontiguousarray(a)


def de_matrix(m, group_id, n_targets, tie_correction=True, use_continuity=True, clip_value=20.0):
    """hpdex_de_matrix on a DeviceMatrix: as de_csc."""
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    n = m.n_cols * int(n_targets)
    outs = [np.empty(max(n, 1)) for _ in range(5)]
    err = ctypes.create_string_buffer(1024)
    rc = lib().hpdex_de_matrix(m.handle, _p(group_id), int(n_targets), int(tie_correction),
                               int(use_continuity), float(clip_value), *[_p(o) for o in outs],
                               err, len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return tuple(o[:n] for o in outs)


def pden_matrix(m, group_id, n_targets, tie_correction=True, continuity_correction=True,
                use_asymptotic=None, use_hist=False, max_bins=100_000, chunk_genes=None,
                clip_value=20.0):
    """hpdex_pden_matrix on a DeviceMatrix: (U_target, P, fc, log2fc, fdr),
    each n_targets x n_cols target-major, flattened."""
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    n = m.n_cols * int(n_targets)
    outs = [np.empty(max(n, 1)) for _ in range(5)]
    o = PdenOpts()
    o.tie_correction = int(bool(tie_correction))
    o.continuity_correction = int(bool(continuity_correction))
    o.use_asymptotic = -1 if use_asymptotic is None else int(bool(use_asymptotic))
    o.use_hist = int(bool(use_hist))
    o.max_bins = int(max_bins)
    o.chunk_genes = int(chunk_genes) if chunk_genes else max(m.n_cols, 1)
    o.clip_value = float(clip_value)
    err = ctypes.create_string_buffer(1024)
    rc = lib().hpdex_pden_matrix(m.handle, _p(group_id), int(n_targets), ctypes.byref(o),
                                 *[_p(x) for x in outs], err, len(err))
    if rc != 0:
        raise HpdexError(err.value.decode())
    return tuple(x[:n] for x in outs)


class DeviceMatrix:
    """A CSC matrix resident in HBM (hpdex_matrix_create), or built on the
    device from CSR / dense input or 32-bit indices (from_scipy, from_dense:
    hpdex_matrix_create_from)."""

    def __init__(self, data, indices, indptr, n_rows):
        data, indices, indptr = _csc_args(data, indices, indptr)
        self.n_rows = int(n_rows)
        self.n_cols = len(indptr) - 1
        self.nnz = len(data)
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_matrix_create(DTYPE_CODES[data.dtype], _p(data), _p(indices), _p(indptr),
                                       self.n_rows, self.n_cols, self.nnz, ctypes.byref(h),
                                       err, len(err))
        if rc != 0:
            raise HpdexError(err.value.decode())
        self._h = h

    @classmethod
    def _create(cls, src, keep):
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_matrix_create_from(ctypes.byref(src), ctypes.byref(h), err, len(err))
        del keep
        if rc != 0:
            raise HpdexError(err.value.decode())
        self._h = h
        R, C, nnz = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        lib().hpdex_matrix_shape(h, ctypes.byref(R), ctypes.byref(C), ctypes.byref(nnz))
        self.n_rows, self.n_cols, self.nnz = int(R.value), int(C.value), int(nnz.value)
        return self

    @classmethod
    def from_arrays(cls, layout, data, indices, indptr, n_rows, n_cols):
        """CSC (indices = rows) or CSR (indices = columns) arrays as they are:
        int32 or int64 indices, any of the ten dtypes; converted on the device."""
        data, indices, indptr = _data_array(data), _index_array(indices), _index_array(indptr)
        if indices.dtype != indptr.dtype:
            indices = indices.astype(np.int64)
            indptr = indptr.astype(np.int64)
        src = MatrixSrc(layout=layout, dtype=DTYPE_CODES[data.dtype], data=_p(data),
                        indices=_p(indices), indptr=_p(indptr),
                        index_bits=64 if indices.dtype == np.int64 else 32,
                        n_rows=int(n_rows), n_cols=int(n_cols), nnz=len(data))
        return cls._create(src, (data, indices, indptr))

    @classmethod
    def from_scipy(cls, m):
        """A scipy CSC or CSR matrix (cells x genes), converted on the device."""
        import scipy.sparse
        if isinstance(m, scipy.sparse.csc_matrix) or getattr(m, "format", None) == "csc":
            layout = LAYOUT_CSC
        elif isinstance(m, scipy.sparse.csr_matrix) or getattr(m, "format", None) == "csr":
            layout = LAYOUT_CSR
        else:
            raise HpdexError("sparse_matrix must be CSR or CSC")
        return cls.from_arrays(layout, m.data, m.indices, m.indptr, m.shape[0], m.shape[1])

    @classmethod
    def from_dense(cls, X):
        """A dense cells x genes array (row-major), its nonzeros kept."""
        X = _data_array(X)
        if X.ndim != 2:
            raise HpdexError("dense input must be 2-D")
        src = MatrixSrc(layout=LAYOUT_DENSE, dtype=DTYPE_CODES[X.dtype], data=_p(X),
                        index_bits=64, n_rows=X.shape[0], n_cols=X.shape[1], nnz=0)
        return cls._create(src, (X,))

    def download(self):
        """(data as float64, rows int32, indptr int64) of the resident CSC."""
        data = np.empty(max(self.nnz, 1))
        rows = np.empty(max(self.nnz, 1), dtype=np.int32)
        indptr = np.empty(self.n_cols + 1, dtype=np.int64)
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_matrix_download(self._h, _p(data), _p(rows), _p(indptr), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        return data[:self.nnz], rows[:self.nnz], indptr

    @property
    def handle(self):
        return self._h

    def device_bytes(self) -> int:
        return int(lib().hpdex_matrix_device_bytes(self._h))

    def ref_cache_info(self) -> dict:
        """The reference tables the matrix keeps: columns with tables and
        their device bytes (both 0 when it keeps none)."""
        cols, nbytes = ctypes.c_int64(), ctypes.c_size_t()
        lib().hpdex_matrix_ref_cache_info(self._h, ctypes.byref(cols), ctypes.byref(nbytes))
        return {"columns": int(cols.value), "bytes": int(nbytes.value)}

    def ref_cache_clear(self):
        lib().hpdex_matrix_ref_cache_clear(self._h)

    def compact_upload(self):
        """(whether the entries crossed PCIe in the 4-byte compact form, the
        columns written as they were because an entry did not qualify)."""
        n = ctypes.c_int64()
        form = lib().hpdex_matrix_upload_form(self._h, ctypes.byref(n))
        return bool(form), int(n.value)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hpdex_matrix_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PLAN_PLAIN, PLAN_BATCHED, PLAN_STATS = 0, 1, 2


class Plan:
    """A job on a DeviceMatrix (hpdex_plan_create); form 0 plain, 1 batched.
    stats=True keeps the tie term and sample sizes for results(with_stats=True)."""

    def __init__(self, matrix: DeviceMatrix, group_id, n_targets, opts: MwuOpts, form=1,
                 stats=False):
        self.matrix = matrix
        self.n_targets = int(n_targets)
        self.stats = bool(stats)
        form = int(form) | (PLAN_STATS if stats else 0)
        gid = np.ascontiguousarray(group_id, dtype=np.int32)
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_plan_create(matrix.handle, _p(gid), self.n_targets, ctypes.byref(opts),
                                     int(form), ctypes.byref(h), err, len(err))
        if rc != 0:
            raise HpdexError(err.value.decode())
        self._h = h

    def run(self, iters=1):
        tot, ker = ctypes.c_double(), ctypes.c_double()
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_plan_run(self._h, int(iters), ctypes.byref(tot), ctypes.byref(ker),
                                  err, len(err))
        if rc != 0:
            raise HpdexError(err.value.decode())
        return tot.value, ker.value

    def results(self, with_stats=False):
        n = self.matrix.n_cols * self.n_targets
        outs = [np.empty(max(n, 1)) for _ in range(5 if with_stats else 2)]
        ptrs = [_p(o) for o in outs] + [None] * (5 - len(outs))
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_plan_results(self._h, *ptrs, err, len(err))
        if rc != 0:
            raise HpdexError(err.value.decode())
        return tuple(o[:n] for o in outs)

    def info(self) -> dict:
        """Column routing of the plan and hand-overs of its last run."""
        i = PlanInfo()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_plan_get_info(self._h, ctypes.byref(i), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        return {n: int(getattr(i, n)) for n, _ in PlanInfo._fields_}

    def stage_times(self, iters=5) -> dict:
        """Mean ms of the count, split, reference and target kernels (HIP
        events between the kernels of the plain launch sequence)."""
        ms = (ctypes.c_double * 4)()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_plan_stage_times(self._h, int(iters), ms, err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        return dict(zip(("count", "split", "reference", "target"), list(ms)))

    KERNELS = ("count", "split", "reference", "target")

    def kernel_ms(self, kernel, iters=20) -> float:
        """Mean ms of one kernel ("count", "split", "reference" or "target")
        over `iters` launches of it back to back between two HIP events."""
        ms = ctypes.c_double()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_plan_kernel_ms(self._h, self.KERNELS.index(kernel), int(iters),
                                      ctypes.byref(ms), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        return ms.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hpdex_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several plans run as one launch sequence (hpdex_batch_*): one queue of
    every plan's count-kernel items, largest first, then the sort-path
    kernels over every plan's other columns, replayed from one hipGraph.
    Results stay in each plan (Plan.results)."""

    def __init__(self, plans):
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_batch_create(arr, len(self.plans), ctypes.byref(h), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        self._h = h

    def run(self, iters=1) -> float:
        """Replay `iters` times and wait; returns the elapsed milliseconds."""
        ms = ctypes.c_double()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_batch_run(self._h, int(iters), ctypes.byref(ms), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        return ms.value

    def info(self) -> dict:
        a, b = ctypes.c_int64(), ctypes.c_int64()
        lib().hpdex_batch_get_info(self._h, ctypes.byref(a), ctypes.byref(b))
        return {"count_items": a.value, "split_columns": b.value}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hpdex_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_tuning() -> Tuning:
    t = Tuning()
    lib().hpdex_get_tuning(ctypes.byref(t))
    return t


def set_tuning(hist_max_bins=None, size_ordered=None, count_lds_bytes=None, count_block=None,
               count_packed=None, split_arrival=None, ref_lds_keys=None, ref_cache=None):
    t = get_tuning()
    if ref_cache is not None:
        t.ref_cache = int(ref_cache)
    if ref_lds_keys is not None:
        t.ref_lds_keys = int(ref_lds_keys)
    if split_arrival is not None:
        t.split_arrival = int(split_arrival)
    if count_packed is not None:
        t.count_packed = int(count_packed)
    if count_block is not None:
        t.count_block = count_block
    if hist_max_bins is not None:
        t.hist_max_bins = hist_max_bins
    if size_ordered is not None:
        t.size_ordered = size_ordered
    if count_lds_bytes is not None:
        t.count_lds_bytes = count_lds_bytes
    err = ctypes.create_string_buffer(1024)
    if lib().hpdex_set_tuning(ctypes.byref(t), err, len(err)) != 0:
        raise HpdexError(err.value.decode())


def device_info():
    cus, lds = ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(256)
    if lib().hpdex_device_info(ctypes.byref(cus), ctypes.byref(lds), name, len(name)) != 0:
        raise HpdexError("no HIP device available")
    return {"cus": cus.value, "max_lds_per_block": lds.value, "name": name.value.decode()}


def bandwidth_probe(nbytes=1 << 30, iters=10) -> float:
    """GB/s of a streaming copy (read + write bytes) on the device."""
    g = ctypes.c_double()
    err = ctypes.create_string_buffer(1024)
    if lib().hpdex_bandwidth_probe(int(nbytes), int(iters), ctypes.byref(g), err, len(err)) != 0:
        raise HpdexError(err.value.decode())
    return g.value


class Session:
    """A serving session (hpdex_session_*): jobs admitted by HBM footprint and
    run concurrently, `slots` at a time, within an arena of `budget_bytes`."""

    def __init__(self, budget_bytes, slots=4):
        o = SessionOpts(int(budget_bytes), int(slots))
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_session_create(ctypes.byref(o), ctypes.byref(h), err, len(err)) != 0:
            raise HpdexError(err.value.decode())
        self._h = h
        self._keep = {}

    def submit(self, group_id, n_targets, opts: MwuOpts, matrix: DeviceMatrix = None,
               data=None, indices=None, indptr=None, n_rows=None, n_cols=None,
               layout=LAYOUT_CSC) -> int:
        """Queue a job on a resident matrix, or on a host CSC (or CSR, with
        n_cols) uploaded for it; int32 indices are sent as they are.  Returns
        the ticket; the arrays stay referenced until wait()."""
        gid = np.ascontiguousarray(group_id, dtype=np.int32)
        j = Job()
        keep = [gid]
        if matrix is not None:
            j.matrix = matrix.handle
            C = matrix.n_cols
            keep.append(matrix)
        else:
            data, indices, indptr = _data_array(data), _index_array(indices), _index_array(indptr)
            if indices.dtype != indptr.dtype:
                indices, indptr = indices.astype(np.int64), indptr.astype(np.int64)
            keep += [data, indices, indptr]
            j.dtype = DTYPE_CODES[data.dtype]
            j.data, j.indices, j.indptr = _p(data), _p(indices), _p(indptr)
            j.index_bits = 32 if indices.dtype == np.int32 else 64
            j.layout = layout
            j.n_rows = len(gid) if n_rows is None else int(n_rows)
            C = len(indptr) - 1 if layout == LAYOUT_CSC else int(n_cols)
            j.n_cols, j.nnz = C, len(data)
        n = C * int(n_targets)
        U1, P = np.empty(max(n, 1)), np.empty(max(n, 1))
        keep += [U1, P]
        j.group_id = _p(gid)
        j.n_targets = int(n_targets)
        j.opts = opts
        j.U1_out, j.P_out = _p(U1), _p(P)
        t = ctypes.c_int64()
        err = ctypes.create_string_buffer(1024)
        if lib().hpdex_session_submit(self._h, ctypes.byref(j), ctypes.byref(t), err,
                                      len(err)) != 0:
            raise HpdexError(err.value.decode())
        self._keep[t.value] = (keep, U1[:n], P[:n])
        return t.value

    def wait(self, ticket):
        """Block until the job ends: (U1, P, times) or HpdexError."""
        tm = JobTimes()
        err = ctypes.create_string_buffer(1024)
        rc = lib().hpdex_session_wait(self._h, int(ticket), ctypes.byref(tm), err, len(err))
        keep, U1, P = self._keep.pop(ticket, (None, None, None))
        if rc != 0:
            raise HpdexError(err.value.decode())
        return U1, P, {n: getattr(tm, n) for n, _ in JobTimes._fields_}

    def peak(self):
        j, b = ctypes.c_int32(), ctypes.c_size_t()
        lib().hpdex_session_peak(self._h, ctypes.byref(j), ctypes.byref(b))
        return {"jobs_in_flight": j.value, "arena_bytes": b.value}

    def batches(self):
        """Merged kernel launches so far, the jobs they ran, the most in one."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        lib().hpdex_session_batches(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"launches": a.value, "jobs": b.value, "max_jobs": c.value}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hpdex_session_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
