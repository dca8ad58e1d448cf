"""CPU oracle for the Mann-Whitney U path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this package, and only as the checker (or the timed CPU baseline).  The product
path (hpdex_amd) never imports it and has no CPU fallback.

The C restatement lives in mwu_oracle.c (it cites the reference lines it
follows); this module builds it on first use with gcc and wraps it with ctypes.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libmwu_oracle.so")
_SRC = os.path.join(_HERE, "mwu_oracle.c")
_lock = threading.Lock()
_lib = None

DTYPE_CODES = {
    np.dtype(np.int8): 0, np.dtype(np.uint8): 1, np.dtype(np.int16): 2,
    np.dtype(np.uint16): 3, np.dtype(np.int32): 4, np.dtype(np.uint32): 5,
    np.dtype(np.int64): 6, np.dtype(np.uint64): 7, np.dtype(np.float32): 8,
    np.dtype(np.float64): 9,
}
ZERO_HANDLING = {"none": 0, "min": 1, "max": 2, "mix": 3}
ALTERNATIVE = {"less": 0, "greater": 1, "two-sided": 2}
METHOD = {"exact": 1, "asymptotic": 2}


def build(force: bool = False) -> str:
    """Compile the oracle with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(_SRC):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _SO


def lib():
    global _lib
    with _lock:
        if _lib is None:
            build()
            L = ctypes.CDLL(_SO)
            dp = ctypes.POINTER(ctypes.c_double)
            vp = ctypes.c_void_p
            sz = ctypes.c_size_t
            ci = ctypes.c_int
            L.oracle_mwu.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz, ci, ci, ci, ci, ci, ci,
                                     vp, vp, vp, vp, vp, ctypes.c_char_p, sz]
            L.oracle_mwu.restype = ci
            L.oracle_group_mean.argtypes = [ci, vp, vp, vp, sz, sz, sz, vp, sz, ci, ci, ci, vp,
                                            ctypes.c_char_p, sz]
            L.oracle_group_mean.restype = ci
            L.oracle_exact_counts.argtypes = [sz, sz, vp, sz]
            L.oracle_exact_counts.restype = ci
            L.oracle_p_asymptotic.argtypes = [ctypes.c_double] * 5 + [ci]
            L.oracle_p_asymptotic.restype = ctypes.c_double
            L.oracle_p_exact.argtypes = [ctypes.c_double, sz, sz, vp, ci]
            L.oracle_p_exact.restype = ctypes.c_double
            del dp
            _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data if a is not None else None


def mwu(data, indices, indptr, group_id, n_targets, n_rows=None, *, tie_correction=True,
        use_continuity=False, zero_handling="mix", alternative="two-sided",
        method="asymptotic", threads=-1, with_stats=False):
    """Reference-algorithm MWU on a CSC triple.  Returns (U1, P) or, with
    with_stats, (U1, P, tie, n1, n2); raises OracleError with the reference's
    message on failure."""
    data = np.ascontiguousarray(data)
    if data.dtype not in DTYPE_CODES:
        raise OracleError("Unsupported dtype for data")
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    R = len(group_id) if n_rows is None else n_rows
    C = len(indptr) - 1
    Np = C * int(n_targets) if n_targets > 0 else 0
    U1 = np.zeros(max(Np, 1)); P = np.ones(max(Np, 1))
    tie = np.zeros(max(Np, 1)); n1 = np.zeros(max(Np, 1)); n2 = np.zeros(max(Np, 1))
    err = ctypes.create_string_buffer(512)
    if threads is None or threads < 0:
        threads = os.cpu_count() or 1
    zh = ZERO_HANDLING[zero_handling] if isinstance(zero_handling, str) else int(zero_handling)
    alt = ALTERNATIVE[alternative] if isinstance(alternative, str) else int(alternative)
    me = METHOD[method] if isinstance(method, str) else int(method)
    rc = lib().oracle_mwu(DTYPE_CODES[data.dtype], _ptr(data), _ptr(indices), _ptr(indptr),
                          R, C, len(data), _ptr(group_id), int(n_targets), int(tie_correction),
                          int(use_continuity), zh, alt, me, int(threads),
                          _ptr(U1), _ptr(P), _ptr(tie), _ptr(n1), _ptr(n2), err, len(err))
    if rc != 0:
        raise OracleError(err.value.decode())
    if with_stats:
        return U1[:Np], P[:Np], tie[:Np], n1[:Np], n2[:Np]
    return U1[:Np], P[:Np]


def group_mean(data, indices, indptr, group_id, n_groups, *, include_zeros=True,
               use_kahan=False, threads=-1):
    data = np.ascontiguousarray(data)
    if data.dtype not in DTYPE_CODES:
        raise OracleError("Unsupported dtype for data")
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    group_id = np.ascontiguousarray(group_id, dtype=np.int32)
    C = len(indptr) - 1
    out = np.zeros(max(C * int(n_groups), 1))
    err = ctypes.create_string_buffer(512)
    if threads is None or threads < 0:
        threads = os.cpu_count() or 1
    rc = lib().oracle_group_mean(DTYPE_CODES[data.dtype], _ptr(data), _ptr(indices), _ptr(indptr),
                                 len(group_id), C, len(data), _ptr(group_id), int(n_groups),
                                 int(include_zeros), int(use_kahan), int(threads), _ptr(out),
                                 err, len(err))
    if rc != 0:
        raise OracleError(err.value.decode())
    return out[:C * int(n_groups)]


def exact_counts(n1: int, n2: int) -> np.ndarray:
    """FP64 counts of the Mann-Whitney null, index = U, length n1*n2+1."""
    out = np.zeros(n1 * n2 + 1)
    if lib().oracle_exact_counts(n1, n2, _ptr(out), 1 << 28) != 0:
        raise OracleError("exact null table too large")
    return out


def p_exact(U1: float, n1: int, n2: int, alternative="two-sided") -> float:
    c = exact_counts(n1, n2)
    return lib().oracle_p_exact(float(U1), n1, n2, _ptr(c), ALTERNATIVE[alternative])


def p_asymptotic(U1, n1, n2, tie, cc, alternative="two-sided") -> float:
    return lib().oracle_p_asymptotic(float(U1), float(n1), float(n2), float(tie), float(cc),
                                     ALTERNATIVE[alternative])
