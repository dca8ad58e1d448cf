"""Python backend wrapper: same names, defaults, validation and messages as the
reference's src/hpdex/backend/mannwhitneyu_.py:1-127, over the MI355X
extension module ``hpdex_amd.backend.bin.kernel``.

Outputs are 1-D read-only float64 arrays of length C*n_targets, gene-major
(index c*n_targets + g-1), exactly as the reference returns them.

Portions reproduce the reference's text (its interface): Copyright (c) 2025
Wang Zhongqi, MIT License -- the full notice is in the repository's NOTICE file.
"""
from typing import Literal

import numpy as np
import scipy.sparse

from .bin import kernel

__all__ = ["mannwhitneyu", "group_mean"]

# ---------------- parameter maps (mannwhitneyu_.py:10-25) ----------------
_zero_handling_map = {
    "none": 0,
    "min": 1,
    "max": 2,
    "mix": 3,
}
_alternative_map = {
    "less": 0,
    "greater": 1,
    "two-sided": 2,
}
_method_map = {
    "exact": 1,
    "asymptotic": 2,
}


def _rows_of(sparse_matrix) -> int:
    # CSC: group_id labels rows; CSR: group_id labels columns (mannwhitneyu_.py:47-51)
    if isinstance(sparse_matrix, scipy.sparse.csc_matrix):
        return sparse_matrix.shape[0]
    return sparse_matrix.shape[1]


def mannwhitneyu(
    sparse_matrix: "scipy.sparse.csr_matrix | scipy.sparse.csc_matrix",
    group_id: np.ndarray,
    n_targets: int,
    ref_sorted: bool = False,
    tar_sorted: bool = False,
    tie_correction: bool = True,
    use_continuity: bool = False,
    fast_norm: bool = False,
    zero_handling: Literal["none", "min", "max", "mix"] = "mix",
    alternative: Literal["less", "greater", "two-sided"] = "two-sided",
    method: Literal["exact", "asymptotic"] = "asymptotic",
    threads: int = -1,
    show_progress: bool = False,
) -> "tuple[np.ndarray, np.ndarray]":
    """Mann-Whitney U test on sparse CSR/CSC matrix (multi-group).

    Reference: src/hpdex/backend/mannwhitneyu_.py:28-86."""
    if not isinstance(sparse_matrix, (scipy.sparse.csr_matrix, scipy.sparse.csc_matrix)):
        raise ValueError("sparse_matrix must be CSR or CSC")
    R = _rows_of(sparse_matrix)
    if group_id.ndim != 1 or len(group_id) != R:
        raise ValueError(
            f"group_id must be 1D with length {R} (rows of CSC / cols of CSR)"
        )
    group_id = group_id.astype(np.int32, copy=False)
    data = sparse_matrix.data
    indices = sparse_matrix.indices.astype(np.int64, copy=False)
    indptr = sparse_matrix.indptr.astype(np.int64, copy=False)

    zero_handling = _zero_handling_map[zero_handling]
    alternative = _alternative_map[alternative]
    method = _method_map[method]

    return kernel.mannwhitneyu(
        data=data,
        indices=indices,
        indptr=indptr,
        group_id=group_id,
        n_targets=n_targets,
        ref_sorted=ref_sorted,
        tar_sorted=tar_sorted,
        tie_correction=tie_correction,
        use_continuity=use_continuity,
        fast_norm=fast_norm,
        zero_handling=zero_handling,
        alternative=alternative,
        method=method,
        threads=threads,
        layout="csc" if isinstance(sparse_matrix, scipy.sparse.csc_matrix) else "csr",
        show_progress=show_progress,
    )


def group_mean(
    sparse_matrix: "scipy.sparse.csr_matrix | scipy.sparse.csc_matrix",
    group_id: np.ndarray,
    n_groups: int,
    include_zeros: bool = True,
    use_kahan: bool = False,
    threads: int = -1,
) -> np.ndarray:
    """Compute group-wise mean for each feature (sparse CSR/CSC).

    Reference: src/hpdex/backend/mannwhitneyu_.py:89-127.  Output index is
    c*n_groups + g."""
    if not isinstance(sparse_matrix, (scipy.sparse.csr_matrix, scipy.sparse.csc_matrix)):
        raise ValueError("sparse_matrix must be CSR or CSC")
    R = _rows_of(sparse_matrix)
    if group_id.ndim != 1 or len(group_id) != R:
        raise ValueError(
            f"group_id must be 1D with length {R} (rows of CSC / cols of CSR)"
        )
    group_id = group_id.astype(np.int32, copy=False)
    data = sparse_matrix.data
    indices = sparse_matrix.indices.astype(np.int64, copy=False)
    indptr = sparse_matrix.indptr.astype(np.int64, copy=False)

    return kernel.group_mean(
        data=data,
        indices=indices,
        indptr=indptr,
        group_id=group_id,
        n_groups=n_groups,
        include_zeros=include_zeros,
        threads=threads,
        use_kahan=use_kahan,
        layout="csc" if isinstance(sparse_matrix, scipy.sparse.csc_matrix) else "csr",
    )
